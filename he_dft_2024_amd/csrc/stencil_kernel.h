// stencil_kernel.h — 13-point finite-difference pencil kernel (derivnD on gfx950).
//
// One workgroup stages whole pencils of the derivative axis in LDS (x rows, or
// tiles of TX consecutive x columns along y or z, loaded with coalesced
// x-contiguous reads), applies the Bickley stencil of derivnD
// (init_deriv_parallel.f90:373-648 real, :933-1209 complex) with the boundary
// rows that redef.f builds for the chosen icon, and hands each derivative value
// to an epilogue functor that stores it or fuses the consumer's arithmetic.
// Input functors fuse the producer's arithmetic into the LDS fill.
#pragma once
#include <cstdlib>
#include <type_traits>
#include <utility>

#include "common.h"

namespace hedft {

// V: element type staged in LDS (double, double2, or float, float2 on the mixed
// path); A = acc_t<V>: the derivative is always accumulated in double.
template <typename V>
struct acc_of {
  using type = V;
};
template <>
struct acc_of<float> {
  using type = double;
};
template <>
struct acc_of<float2> {
  using type = double2;
};
template <typename V>
using acc_t = typename acc_of<V>::type;

template <typename V>
__device__ __forceinline__ V vzero();
template <>
__device__ __forceinline__ double vzero<double>() { return 0.0; }
template <>
__device__ __forceinline__ double2 vzero<double2>() { return make_double2(0.0, 0.0); }
template <>
__device__ __forceinline__ float vzero<float>() { return 0.0f; }
template <>
__device__ __forceinline__ float2 vzero<float2>() { return make_float2(0.0f, 0.0f); }
__device__ __forceinline__ double vfma(double c, double v, double acc) { return fma(c, v, acc); }
__device__ __forceinline__ double2 vfma(double c, double2 v, double2 acc) {
  return make_double2(fma(c, v.x, acc.x), fma(c, v.y, acc.y));
}
__device__ __forceinline__ double vfma(double c, float v, double acc) { return fma(c, (double)v, acc); }
__device__ __forceinline__ double2 vfma(double c, float2 v, double2 acc) {
  return make_double2(fma(c, (double)v.x, acc.x), fma(c, (double)v.y, acc.y));
}
__device__ __forceinline__ double vmul(double c, double v) { return c * v; }
__device__ __forceinline__ double2 vmul(double c, double2 v) { return make_double2(c * v.x, c * v.y); }
// sign flips of periodic halos (c = +-1, exact)
__device__ __forceinline__ float vmul(double c, float v) { return (float)c * v; }
__device__ __forceinline__ float2 vmul(double c, float2 v) { return make_float2((float)c * v.x, (float)c * v.y); }

// Derivative at position i of a pencil P (element stride s) of length n.
template <typename V>
__device__ __forceinline__ acc_t<V> stencil_at(const StencilCoef& cf, const V* P, int s, int i, int n) {
  acc_t<V> acc = vzero<acc_t<V>>();
  if (cf.mode != 0) {
    const double sg = (cf.mode == 2) ? -1.0 : 1.0;
#pragma unroll
    for (int j = 0; j < 13; ++j) {
      int ii = i + j - 6;
      double c = cf.c[j];
      if (ii < 0) { ii += n; c *= sg; }
      else if (ii >= n) { ii -= n; c *= sg; }
      acc = vfma(c, P[ii * s], acc);
    }
  } else if (i < 6) {
#pragma unroll
    for (int j = 0; j < 13; ++j) acc = vfma(cf.rows[i][j], P[j * s], acc);
  } else if (i >= n - 6) {
    const int r = i - n + 12;
#pragma unroll
    for (int j = 0; j < 13; ++j) acc = vfma(cf.rows[r][j], P[(n - 13 + j) * s], acc);
  } else {
#pragma unroll
    for (int j = 0; j < 13; ++j) acc = vfma(cf.c[j], P[(i - 6 + j) * s], acc);
  }
  return acc;
}

// Sliding-window evaluation of KB consecutive outputs i0..i0+KB-1 of a pencil
// (element stride s, length n): the 12+KB inputs are read once into registers.
// Outputs whose 13-point window needs a boundary row fall back to stencil_at.
template <typename V, int KB>
__device__ __forceinline__ void stencil_run(const StencilCoef& cf, const V* P, int s, int i0, int cnt, int n,
                                            acc_t<V> (&out)[KB]) {
  const bool interior = (cf.mode != 0) || (i0 >= 6 && i0 + KB <= n - 6);
  if (interior && cnt == KB) {
    V w[KB + 12];
    const double sg = (cf.mode == 2) ? -1.0 : 1.0;
#pragma unroll
    for (int q = 0; q < KB + 12; ++q) {
      int ii = i0 - 6 + q;
      double f = 1.0;
      if (ii < 0) { ii += n; f = sg; }
      else if (ii >= n) { ii -= n; f = sg; }
      w[q] = vmul(f, P[ii * s]);
    }
#pragma unroll
    for (int k = 0; k < KB; ++k) {
      acc_t<V> acc = vzero<acc_t<V>>();
#pragma unroll
      for (int j = 0; j < 13; ++j) acc = vfma(cf.c[j], w[k + j], acc);
      out[k] = acc;
    }
  } else {
#pragma unroll
    for (int k = 0; k < KB; ++k)
      if (k < cnt) out[k] = stencil_at<V>(cf, P, s, i0 + k, n);
  }
}

// Outputs per sliding window: Out::KB where the epilogue declares it (heavy
// epilogues keep fewer outputs in flight to stay within registers), else 8.
template <class Out, class = void>
struct out_kb {
  static constexpr int value = 8;
};
template <class Out>
struct out_kb<Out, std::void_t<decltype(Out::KB)>> {
  static constexpr int value = Out::KB;
};

// Epilogues that declare CENTRE = true receive the pencil's value at the
// output point (from the LDS tile) with the derivative.
template <class Out, class = void>
struct out_centre {
  static constexpr bool value = false;
};
template <class Out>
struct out_centre<Out, std::void_t<decltype(Out::CENTRE)>> {
  static constexpr bool value = Out::CENTRE;
};

// Generic pencil kernel.  AX: 0 = x, 1 = y, 2 = z.  In::load(idx) -> V,
// Out::store(idx, ix, iy, iz, d, red) with Out::NR per-thread partial sums
// reduced per workgroup into Out::partials.  For AX = 1, 2 a workgroup owns TX
// consecutive x columns of whole pencils; each thread evaluates a contiguous
// run of outputs along the pencil with a register window.
template <typename V, int AX, int TX, class In, class Out>
__global__ void __launch_bounds__(256) k_pencil(In in, Out out, StencilCoef cf, Dims d, int tx) {
  extern __shared__ unsigned char smem_raw[];
  V* lds = reinterpret_cast<V*>(smem_raw);
  constexpr int NR = Out::NR;
  double red[NR > 0 ? NR : 1];
#pragma unroll
  for (int r = 0; r < (NR > 0 ? NR : 1); ++r) red[r] = 0.0;
  const int n = (AX == 0) ? d.nx : (AX == 1) ? d.ny : d.nz;
  if (AX == 0) {
    const int rows = tx;
    const int row0 = blockIdx.x * rows;
    const int nrows = d.ny * d.nz;
    for (int e = threadIdx.x; e < rows * n; e += blockDim.x) {
      const int r = e / n, i = e - r * n;
      const int row = row0 + r;
      lds[e] = (row < nrows) ? in.load((size_t)row * n + i) : vzero<V>();
    }
    __syncthreads();
    for (int e = threadIdx.x; e < rows * n; e += blockDim.x) {
      const int r = e / n, i = e - r * n;
      const int row = row0 + r;
      if (row < nrows) {
        const acc_t<V> v = stencil_at<V>(cf, lds + r * n, 1, i, n);
        if constexpr (out_centre<Out>::value)
          out.store((size_t)row * n + i, i, row % d.ny, row / d.ny, v, lds[e], red);
        else
          out.store((size_t)row * n + i, i, row % d.ny, row / d.ny, v, red);
      }
    }
  } else {
    constexpr int TPC = 256 / TX, KB = out_kb<Out>::value;
    const int x0 = blockIdx.x * TX;
    size_t base, stride;
    if (AX == 1) {
      base = (size_t)blockIdx.y * d.nx * d.ny + x0;
      stride = d.nx;
    } else {
      base = (size_t)blockIdx.y * d.nx + x0;
      stride = (size_t)d.nx * d.ny;
    }
    const int xl = threadIdx.x % TX, t = threadIdx.x / TX;
    const bool col_ok = x0 + xl < d.nx;
    // element e = j*TX + xl of the tile; thread order keeps x fastest (coalesced)
    const int cmax = d.nx - 1 - x0;  // columns past the grid edge are loaded clamped and never stored
    lds_fill<8, 256>(lds, n * TX, [&](int e) {
      const int j = e / TX, c = min(e % TX, cmax);
      return in.load(base + j * stride + c);
    });
    __syncthreads();
    if (col_ok) {
      const int run = n / TPC, i_beg = t * run;
      for (int i0 = i_beg; i0 < i_beg + run; i0 += KB) {
        const int cnt = (i_beg + run - i0 < KB) ? i_beg + run - i0 : KB;
        acc_t<V> v[KB];
        stencil_run<V, KB>(cf, lds + xl, TX, i0, cnt, n, v);
#pragma unroll
        for (int k = 0; k < KB; ++k) {
          if (k < cnt) {
            const int i = i0 + k;
            const int iy = (AX == 1) ? i : (int)blockIdx.y;
            const int iz = (AX == 1) ? (int)blockIdx.y : i;
            if constexpr (out_centre<Out>::value)
              out.store(base + i * stride + xl, x0 + xl, iy, iz, v[k], lds[i * TX + xl], red);
            else
              out.store(base + i * stride + xl, x0 + xl, iy, iz, v[k], red);
          }
        }
      }
    }
  }
  if constexpr (NR > 0) block_reduce_store<NR>(red, out.partials);
}

struct PencilGeom {
  dim3 grid, block;
  size_t lds;
  int tx;
  size_t nblocks;
};

// side-stream stencil passes sized to share CUs with the plane FFTs, on grids
// whose square planes take the whole-plane kernels (HEDFT_SIDE_FIT=0 in the
// environment selects the wider tiles everywhere)
inline bool side_fit(const Dims& d) {
  static const int enabled = [] {
    const char* e = std::getenv("HEDFT_SIDE_FIT");
    return (e && e[0] == '0') ? 0 : 1;
  }();
  return enabled && d.nx == d.ny && d.nx <= 128;
}

// column-tile width of the y/z pencil kernels: 32 float columns (128 B rows;
// 64 columns left a 128^3 grid only 256 workgroups, measured 4 % slower on
// the mixed step), 32 double, 16 double2 (64 KB of LDS at n = 256, where 8
// columns measured 5-7 % slower on the 256^3 steps).  Double columns are 16
// wide (128 B rows) when side passes fit beside the plane FFTs: 16 KiB of LDS
// at n = 128 instead of 32 KiB, under the 30 KiB a plane workgroup leaves free.
template <typename V>
inline int pencil_tx(const Dims& d, int n) {
  if (sizeof(V) == 4) return 32;
  if (sizeof(V) == 8) return side_fit(d) ? 16 : 32;
  return 16;
}

template <typename V>
inline PencilGeom pencil_geom(const Dims& d, int ax) {
  PencilGeom g;
  g.block = dim3(256);
  if (ax == 0) {
    int rows = 256 / d.nx;
    if (rows < 1) rows = 1;
    g.tx = rows;
    const int nrows = d.ny * d.nz;
    g.grid = dim3((nrows + rows - 1) / rows);
    g.lds = sizeof(V) * (size_t)rows * d.nx;
  } else {
    const int n = (ax == 1) ? d.ny : d.nz;
    const int tx = pencil_tx<V>(d, n);
    g.tx = tx;
    g.grid = dim3((d.nx + tx - 1) / tx, (ax == 1) ? d.nz : d.ny);
    g.lds = sizeof(V) * (size_t)tx * n;
  }
  g.nblocks = (size_t)g.grid.x * g.grid.y * g.grid.z;
  return g;
}

template <typename V, int AX, class In, class Out>
inline void launch_pencil(hipStream_t s, const Dims& d, const StencilCoef& cf, const In& in,
                          const Out& out) {
  PencilGeom g = pencil_geom<V>(d, AX);
  if (AX != 0 && g.tx == 64)
    hipLaunchKernelGGL((k_pencil<V, AX, 64, In, Out>), g.grid, g.block, g.lds, s, in, out, cf, d, g.tx);
  else if (AX == 0 || g.tx == 32)
    hipLaunchKernelGGL((k_pencil<V, AX, 32, In, Out>), g.grid, g.block, g.lds, s, in, out, cf, d, g.tx);
  else if (g.tx == 16)
    hipLaunchKernelGGL((k_pencil<V, AX, 16, In, Out>), g.grid, g.block, g.lds, s, in, out, cf, d, g.tx);
  else
    hipLaunchKernelGGL((k_pencil<V, AX, 8, In, Out>), g.grid, g.block, g.lds, s, in, out, cf, d, g.tx);
}

// ---- fused x + y derivatives ------------------------------------------------
// A workgroup holds a tile of one z-plane: tx columns plus a 6-point halo on
// each side in x, all ny rows.  The x derivative reads its 13-point window along
// the tile row, the y derivative the whole column; boundary rows follow the
// same rules as stencil_at (periodic halos are loaded wrapped, with the sign
// flip of antiperiodic icon=18 applied to the wrapped values).
template <typename V>
__device__ __forceinline__ acc_t<V> stencil_tile_x(const StencilCoef& cf, const V* row, int gx, int nx, int x0) {
  acc_t<V> acc = vzero<acc_t<V>>();
  const int off = 6 - x0;  // LDS position of global index g is g + off
  if (cf.mode != 0 || (gx >= 6 && gx < nx - 6)) {
#pragma unroll
    for (int j = 0; j < 13; ++j) acc = vfma(cf.c[j], row[gx - 6 + j + off], acc);
  } else if (gx < 6) {
#pragma unroll
    for (int j = 0; j < 13; ++j) acc = vfma(cf.rows[gx][j], row[j + off], acc);
  } else {
    const int r = gx - nx + 12;
#pragma unroll
    for (int j = 0; j < 13; ++j) acc = vfma(cf.rows[r][j], row[nx - 13 + j + off], acc);
  }
  return acc;
}

// In::loadx(i) feeds the x derivative, In::loady(i) the y derivative (SAME:
// the same operand, staged once)
template <typename V, bool SAME, int TX, class In, class Out>
__global__ void __launch_bounds__(256) k_pencil_xy(In in, Out out, StencilCoef cfx, StencilCoef cfy, Dims d,
                                                   int tx_unused) {
  extern __shared__ unsigned char smem_raw[];
  constexpr int NR = Out::NR;
  double red[NR > 0 ? NR : 1];
#pragma unroll
  for (int r = 0; r < (NR > 0 ? NR : 1); ++r) red[r] = 0.0;
  constexpr int TW = TX + 12, TPC = 256 / TX, KB = out_kb<Out>::value;
  const int nx = d.nx, ny = d.ny;
  V* A = reinterpret_cast<V*>(smem_raw);
  V* B = SAME ? A : A + TW * ny;
  // XCD-aware tile order: workgroups are dispatched round-robin over the 8
  // XCDs, so physical id p runs on XCD p % 8.  Giving each XCD a contiguous
  // range of tiles (x tiles of a plane next to each other) lets the 12 halo
  // columns a tile shares with its x neighbours hit that XCD's L2.
  const int ntiles = gridDim.x * gridDim.y;
  int tile = blockIdx.x + gridDim.x * blockIdx.y;
  if ((ntiles & 7) == 0) tile = (tile & 7) * (ntiles >> 3) + (tile >> 3);
  const int x0 = (tile % gridDim.x) * TX, z = tile / gridDim.x;
  const size_t zb = (size_t)z * nx * ny;
  {
    // Halo slots outside the box (non-periodic modes) are never read by any
    // output, so they are filled from a clamped in-range address; periodic
    // halos hold the wrapped value times the antiperiodic sign.  The y operand
    // (B, when it differs from the x operand) is only needed on the TX centre
    // columns.
    constexpr int U = 8;
    const double sgw = (cfx.mode == 2) ? -1.0 : 1.0;
    for (int e0 = threadIdx.x; e0 < TW * ny; e0 += 256 * U) {
      V ta[U];
      double sg[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = min(e0 + u * 256, TW * ny - 1);
        const int y = e / TW, xl = e - y * TW;
        int gx = x0 - 6 + xl;
        sg[u] = 1.0;
        if (gx < 0 || gx >= nx) {
          if (cfx.mode != 0) {
            // one wrap covers every slot an output reads; a tile wider than
            // the grid (tx > nx) has slots beyond that, which are clamped so
            // that their (unused) loads stay inside the array
            sg[u] = sgw;
            gx = min(max(gx + ((gx < 0) ? nx : -nx), 0), nx - 1);
          } else {
            gx = min(max(gx, 0), nx - 1);
          }
        }
        ta[u] = in.loadx(zb + (size_t)y * nx + gx);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = e0 + u * 256;
        if (e < TW * ny) A[e] = vmul(sg[u], ta[u]);
      }
    }
    if constexpr (!SAME) {
      for (int e0 = threadIdx.x; e0 < TX * ny; e0 += 256 * U) {
        V tb[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int e = min(e0 + u * 256, TX * ny - 1);
          const int y = e / TX, xl = e - y * TX;
          tb[u] = in.loady(zb + (size_t)y * nx + min(x0 + xl, nx - 1));
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int e = e0 + u * 256;
          if (e < TX * ny) {
            const int y = e / TX, xl = e - y * TX;
            B[y * TW + xl + 6] = tb[u];
          }
        }
      }
    }
  }
  __syncthreads();
  const int xl = threadIdx.x % TX, t = threadIdx.x / TX;
  const int gx = x0 + xl;
  if (gx < nx) {
    const int run = ny / TPC, i_beg = t * run;
    for (int y0 = i_beg; y0 < i_beg + run; y0 += KB) {
      const int cnt = (i_beg + run - y0 < KB) ? i_beg + run - y0 : KB;
      acc_t<V> dy[KB];
      stencil_run<V, KB>(cfy, B + xl + 6, TW, y0, cnt, ny, dy);
#pragma unroll
      for (int k = 0; k < KB; ++k) {
        if (k < cnt) {
          const int y = y0 + k;
          const acc_t<V> dx = stencil_tile_x<V>(cfx, A + y * TW, gx, nx, x0);
          out.store(zb + (size_t)y * nx + gx, gx, y, z, dx, dy[k], red);
        }
      }
    }
  }
  if constexpr (NR > 0) block_reduce_store<NR>(red, out.partials);
}

// 13-point derivative at global index g of an axis of length n, from a tile
// whose local element k holds global index g0 - 6 + k (stride s); halos are
// preloaded (wrapped with the antiperiodic sign), so no index wraps here.
template <typename V>
__device__ __forceinline__ acc_t<V> stencil_tile(const StencilCoef& cf, const V* P, int s, int g, int n, int g0) {
  acc_t<V> acc = vzero<acc_t<V>>();
  const int off = 6 - g0;
  if (cf.mode != 0 || (g >= 6 && g < n - 6)) {
#pragma unroll
    for (int j = 0; j < 13; ++j) acc = vfma(cf.c[j], P[(g - 6 + j + off) * s], acc);
  } else if (g < 6) {
#pragma unroll
    for (int j = 0; j < 13; ++j) acc = vfma(cf.rows[g][j], P[(j + off) * s], acc);
  } else {
    const int r = g - n + 12;
#pragma unroll
    for (int j = 0; j < 13; ++j) acc = vfma(cf.rows[r][j], P[(n - 13 + j + off) * s], acc);
  }
  return acc;
}

// x + y derivatives on 2-D tiles: TX columns x TY rows of one z-plane with
// 6-point halos on all four sides, so that wide tiles (few halo re-reads)
// fit in LDS at any ny.  SAME: one (TY+12) x (TX+12) rectangle serves both
// derivatives; otherwise the x operand is staged on TY rows x (TX+12) and the
// y operand on (TY+12) rows x TX.  Requires ny % TY == 0 (every row chunk then
// holds the 13 rows a one-sided boundary stencil reads) and Out::NR == 0.
template <typename V, bool SAME, int TX, int TY, class In, class Out>
__global__ void __launch_bounds__(256) k_pencil_xy2(In in, Out out, StencilCoef cfx, StencilCoef cfy, Dims d) {
  extern __shared__ unsigned char smem_raw[];
  static_assert(Out::NR == 0, "2-D tiles carry no reductions");
  constexpr int TW = TX + 12, TH = TY + 12, TPC = 256 / TX, RUN = TY / TPC;
  static_assert(TY % TPC == 0, "rows per thread");
  const int nx = d.nx, ny = d.ny;
  V* A = reinterpret_cast<V*>(smem_raw);       // x operand, column 0 <-> x0 - 6
  V* B = SAME ? A + 6 : A + TY * TW;           // y operand, row 0 <-> y0 - 6
  constexpr int SB = SAME ? TW : TX;           // row stride of the y operand
  // XCD-aware tile order (see k_pencil_xy): x tiles, then y tiles, then planes
  const int ntx = gridDim.x, nty = gridDim.y;
  const int ntiles = ntx * nty * gridDim.z;
  int tile = blockIdx.x + ntx * (blockIdx.y + nty * blockIdx.z);
  if ((ntiles & 7) == 0) tile = (tile & 7) * (ntiles >> 3) + (tile >> 3);
  const int x0 = (tile % ntx) * TX, y0 = ((tile / ntx) % nty) * TY, z = tile / (ntx * nty);
  const size_t zb = (size_t)z * nx * ny;
  const double sgx = (cfx.mode == 2) ? -1.0 : 1.0, sgy = (cfy.mode == 2) ? -1.0 : 1.0;
  // global index -> in-range index and sign: periodic halos wrap (with the
  // antiperiodic sign), others are clamped (never read by any output).  One
  // wrap covers every slot an output reads; a tile wider than the grid (TX >
  // nx, e.g. 64 columns on a 32-point axis) has slots beyond that, which are
  // clamped too so that their unused loads stay inside the array.
  auto fold = [](int g, int n, int mode, double sgw, double& sg) {
    if (g < 0 || g >= n) {
      if (mode != 0) {
        sg *= sgw;
        return min(max(g + ((g < 0) ? n : -n), 0), n - 1);
      }
      return min(max(g, 0), n - 1);
    }
    return g;
  };
  {
    constexpr int U = 8;
    // x operand (SAME: the whole rectangle)
    constexpr int RA = SAME ? TH : TY;
    for (int e0 = threadIdx.x; e0 < RA * TW; e0 += 256 * U) {
      V ta[U];
      double sg[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = min(e0 + u * 256, RA * TW - 1);
        const int r = e / TW, c = e - r * TW;
        sg[u] = 1.0;
        const int gy = fold(SAME ? y0 - 6 + r : y0 + r, ny, cfy.mode, sgy, sg[u]);
        const int gx = fold(x0 - 6 + c, nx, cfx.mode, sgx, sg[u]);
        ta[u] = in.loadx(zb + (size_t)gy * nx + gx);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = e0 + u * 256;
        if (e < RA * TW) A[e] = vmul(sg[u], ta[u]);
      }
    }
    if constexpr (!SAME) {
      for (int e0 = threadIdx.x; e0 < TH * TX; e0 += 256 * U) {
        V tb[U];
        double sg[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int e = min(e0 + u * 256, TH * TX - 1);
          const int r = e / TX, c = e - r * TX;
          sg[u] = 1.0;
          const int gy = fold(y0 - 6 + r, ny, cfy.mode, sgy, sg[u]);
          tb[u] = in.loady(zb + (size_t)gy * nx + min(x0 + c, nx - 1));
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int e = e0 + u * 256;
          if (e < TH * TX) B[e] = vmul(sg[u], tb[u]);
        }
      }
    }
  }
  __syncthreads();
  const int xl = threadIdx.x % TX, t = threadIdx.x / TX;
  const int gx = x0 + xl;
  if (gx < nx) {
#pragma unroll 2
    for (int k = 0; k < RUN; ++k) {
      const int y = y0 + t * RUN + k;
      const V* rowx = A + (size_t)(y - y0 + (SAME ? 6 : 0)) * TW;
      const acc_t<V> dx = stencil_tile<V>(cfx, rowx, 1, gx, nx, x0);
      const acc_t<V> dy = stencil_tile<V>(cfy, B + xl, SB, y, ny, y0);
      double* no_red = nullptr;
      out.store(zb + (size_t)y * nx + gx, gx, y, z, dx, dy, no_red);
    }
  }
}

// tile shape of the 2-D variant for a value type: (TX, TY).  FIT selects
// tiles whose LDS (at most 30 KiB) fits beside a whole-plane FFT workgroup
// (130 KiB of the CU's 160 KiB), so that a stencil pass on the side stream
// can share CUs with the plane transforms of the main stream.
template <typename V, bool SAME, bool FIT>
struct XYTile {
  static constexpr int TX = (sizeof(V) == 4) ? 64 : (sizeof(V) == 8) ? ((FIT && !SAME) ? 32 : 64) : 32;
  static constexpr int TY = (sizeof(V) == 4) ? 64 : (sizeof(V) == 8) ? 32 : (FIT ? 16 : 32);
  static constexpr size_t LDS = sizeof(V) * (SAME ? (size_t)(TY + 12) * (TX + 12)
                                                 : (size_t)TY * (TX + 12) + (size_t)(TY + 12) * TX);
  static_assert(!FIT || sizeof(V) == 4 || LDS <= 30 * 1024, "FIT tile exceeds the LDS left beside a plane");
};

// 1 when the 2-D tiles are used for this grid (ny a multiple of TY, and a
// last x tile of at least 7 columns); HEDFT_XY_TILE=0 in the environment
// selects the full-column kernel
template <typename V, bool SAME, bool FIT>
inline bool use_xy2(const Dims& d) {
  static const int enabled = [] {
    const char* e = std::getenv("HEDFT_XY_TILE");
    return (e && e[0] == '0') ? 0 : 1;
  }();
  // the last x tile must hold the 13 columns a one-sided boundary stencil reads
  const int rx = d.nx % XYTile<V, SAME, FIT>::TX;
  return enabled && d.ny % XYTile<V, SAME, FIT>::TY == 0 && d.nx >= 13 && (rx == 0 || rx >= 7);
}

template <typename V, bool SAME>
inline PencilGeom pencil_xy_geom(const Dims& d) {
  PencilGeom g;
  g.block = dim3(256);
  int tx = (sizeof(V) == 4) ? 64 : (sizeof(V) == 8) ? 32 : 16;
  auto lds = [&](int t) { return sizeof(V) * (size_t)(t + 12) * d.ny * (SAME ? 1 : 2); };
  while (tx > 8 && lds(tx) > 64 * 1024) tx /= 2;
  // the last tile must start at or before nx-7 so that the one-sided rows of the
  // last six points read inside the tile (true for every supported edge)
  g.tx = tx;
  g.grid = dim3((d.nx + tx - 1) / tx, d.nz);
  g.lds = lds(tx);
  g.nblocks = (size_t)g.grid.x * g.grid.y;
  return g;
}

template <typename V, bool SAME, int TX, class In, class Out>
inline void launch_pencil_xy_t(hipStream_t s, const PencilGeom& g, const Dims& d, const StencilCoef& cfx,
                               const StencilCoef& cfy, const In& in, const Out& out) {
  static bool attr = false;
  if (!attr && g.lds > 64 * 1024) {
    hipFuncSetAttribute((const void*)k_pencil_xy<V, SAME, TX, In, Out>,
                        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    attr = true;
  }
  hipLaunchKernelGGL((k_pencil_xy<V, SAME, TX, In, Out>), g.grid, g.block, g.lds, s, in, out, cfx, cfy, d, g.tx);
}

template <typename V, bool SAME, class In, class Out>
inline void launch_pencil_xy(hipStream_t s, const Dims& d, const StencilCoef& cfx, const StencilCoef& cfy,
                             const In& in, const Out& out) {
  if constexpr (Out::NR == 0) {
    auto xy2 = [&](auto fit) {
      constexpr bool FIT = decltype(fit)::value;
      using T = XYTile<V, SAME, FIT>;
      const dim3 grid((d.nx + T::TX - 1) / T::TX, d.ny / T::TY, d.nz);
      hipLaunchKernelGGL((k_pencil_xy2<V, SAME, T::TX, T::TY, In, Out>), grid, dim3(256), T::LDS, s, in, out, cfx,
                         cfy, d);
    };
    if (side_fit(d)) {
      if (use_xy2<V, SAME, true>(d)) return xy2(std::true_type{});
    } else if (use_xy2<V, SAME, false>(d)) {
      return xy2(std::false_type{});
    }
  }
  PencilGeom g = pencil_xy_geom<V, SAME>(d);
  if (g.tx == 64) launch_pencil_xy_t<V, SAME, 64>(s, g, d, cfx, cfy, in, out);
  else if (g.tx == 32) launch_pencil_xy_t<V, SAME, 32>(s, g, d, cfx, cfy, in, out);
  else if (g.tx == 16) launch_pencil_xy_t<V, SAME, 16>(s, g, d, cfx, cfy, in, out);
  else launch_pencil_xy_t<V, SAME, 8>(s, g, d, cfx, cfy, in, out);
}

// Out types whose H carries classical impurities (Out::h.im)
template <class Out, class = void>
struct has_imp : std::false_type {};
template <class Out>
struct has_imp<Out, std::void_t<decltype(std::declval<Out&>().h.im)>> : std::true_type {};

// ---- 2.5-D fused Laplacian of a complex field with a z-pass epilogue --------
// A workgroup owns a TX x TY tile of (x, y) columns over the planes z0..z1-1
// (z1 - z0 <= tz).  Each thread keeps the 13 planes of its column that the z
// stencil reads in registers and moves up one plane per step.  Each plane of
// the tile, with 6-point x and y halos (periodic halos wrapped with the
// antiperiodic sign, others clamped: no output reads those), is staged in LDS,
// double-buffered, for the x and y derivatives; the halo and the next queue
// value of plane z+1 are loaded while plane z is evaluated.  The epilogue
// (Out::apply, with the operands Out::load read one step ahead)
// receives (d_x + d_y) + d_z and the value at the point: the same sums, in
// the same order, as the x+y tile pass (k_pencil_xy2 with OutSumXY) followed
// by the z pass (k_pencil) produce.  cf3: the x, y, z coefficient sets in
// device memory.  Requires nx % TX == 0, ny % TY == 0, TX, TY >= 7 (a tile
// at the box edge then holds the 13 rows a one-sided stencil reads) and
// n >= 13 on every axis.
template <int TX, int TY, class Out>
__global__ void __launch_bounds__(TX * TY) k_lap25(const double2* __restrict__ P, Out out,
                                                  const StencilCoef* __restrict__ cf3, Dims d, int tz) {
  static_assert(TX >= 7 && TY >= 7, "edge tiles must hold 13 rows");
  constexpr int NT = TX * TY, TW = TX + 12, TH = TY + 12;
  // halo slots: the 6-row bands above and below the centre columns, and the
  // 6-column bands left and right of the centre rows (the corners are read by
  // no stencil)
  constexpr int NHALO = 12 * TX + 12 * TY, HU = (NHALO + NT - 1) / NT;
  __shared__ double2 buf[2][TH * TW];
  // halo ring: the halos of planes z+1..z+6 at step z, plane p in slot p % 6.
  // A plane's halo is read from global memory when the plane enters the z
  // window (the step its neighbouring tiles read the same plane into theirs,
  // so it comes from the L2) and parked here until the plane is staged; each
  // thread reads back only the entries it wrote
  __shared__ double2 ring[6 * NHALO];
  constexpr int NR = Out::NR;
  double red[NR > 0 ? NR : 1];
#pragma unroll
  for (int r = 0; r < (NR > 0 ? NR : 1); ++r) red[r] = 0.0;
  // the coefficients live in LDS: read from global memory inside the loop
  // they would be vector loads, whose waits would also wait for the prefetches
  __shared__ StencilCoef cfs[3];
  static_assert(sizeof(StencilCoef) % 8 == 0, "copied as doubles");
  for (int e = threadIdx.x; e < (int)(3 * sizeof(StencilCoef) / 8); e += NT)
    reinterpret_cast<double*>(cfs)[e] = reinterpret_cast<const double*>(cf3)[e];
  Out o = out;
  __shared__ double rimp_lds[3 * HEDFT_MAX_IMP];
  if constexpr (has_imp<Out>::value) {
    // impurity positions (fixed during the pass) from LDS: no global load
    // of them inside the loop
    if ((int)threadIdx.x < 3 * o.h.im.nimp) rimp_lds[threadIdx.x] = o.h.im.rimp[threadIdx.x];
    o.h.im.rimp = rimp_lds;
  }
  __syncthreads();
  const StencilCoef& cfx = cfs[0];
  const StencilCoef& cfy = cfs[1];
  const StencilCoef& cfz = cfs[2];
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const int ntx = nx / TX, nty = ny / TY, nzc = (nz + tz - 1) / tz;
  // XCD-aware order (see k_pencil_xy): each XCD takes a contiguous range of
  // tiles, x tiles then y tiles then z chunks, so the halos a tile shares
  // with its x and y neighbours stay in that XCD's L2
  const int ntiles = ntx * nty * nzc;
  int tile = blockIdx.x;
  if ((ntiles & 7) == 0) tile = (tile & 7) * (ntiles >> 3) + (tile >> 3);
  const int x0 = (tile % ntx) * TX, y0 = ((tile / ntx) % nty) * TY, zc = tile / (ntx * nty);
  const int z0 = zc * tz, z1 = min(z0 + tz, nz);
  const int xl = threadIdx.x % TX, yl = threadIdx.x / TX;
  const int gx = x0 + xl, gy = y0 + yl;
  const size_t plane = (size_t)nx * ny;
  const size_t col = (size_t)gy * nx + gx;
  const int cpos = (yl + 6) * TW + xl + 6;
  const double sgx = (cfx.mode == 2) ? -1.0 : 1.0, sgy = (cfy.mode == 2) ? -1.0 : 1.0;
  const double sgz = (cfz.mode == 2) ? -1.0 : 1.0;
  int hoff[HU];
  size_t hsrc[HU];
  double hsg[HU];
#pragma unroll
  for (int u = 0; u < HU; ++u) {
    const int k = min((int)threadIdx.x + u * NT, NHALO - 1);
    int r, c;
    if (k < 6 * TX) {
      r = k / TX;
      c = 6 + k % TX;
    } else if (k < 12 * TX) {
      r = TY + 6 + (k - 6 * TX) / TX;
      c = 6 + (k - 6 * TX) % TX;
    } else {
      const int k3 = k - 12 * TX, cc = k3 % 12;
      r = 6 + k3 / 12;
      c = (cc < 6) ? cc : TX + cc;
    }
    double sg = 1.0;
    int yy = y0 - 6 + r, xx = x0 - 6 + c;
    if (yy < 0 || yy >= ny) {
      if (cfy.mode != 0) {
        sg *= sgy;
        yy += (yy < 0) ? ny : -ny;
      } else {
        yy = min(max(yy, 0), ny - 1);
      }
    }
    if (xx < 0 || xx >= nx) {
      if (cfx.mode != 0) {
        sg *= sgx;
        xx += (xx < 0) ? nx : -nx;
      } else {
        xx = min(max(xx, 0), nx - 1);
      }
    }
    hoff[u] = ((int)threadIdx.x + u * NT < NHALO) ? r * TW + c : -1;
    hsrc[u] = (size_t)yy * nx + xx;
    hsg[u] = sg;
  }
  // plane of queue slot 0 at step z (unwrapped; clamped to the box for the
  // one-sided rows of mode 0)
  const bool clampz = cfz.mode == 0;
  auto qbase = [&](int z) { return clampz ? min(max(z - 6, 0), nz - 13) : z - 6; };
  // unwrapped plane u -> stored plane, and the antiperiodic sign of the wrap
  auto zwrap = [&](int u, double& f) {
    f = 1.0;
    if (u < 0) {
      f = sgz;
      return u + nz;
    }
    if (u >= nz) {
      f = sgz;
      return u - nz;
    }
    return u;
  };
  double2 q[13];
  int b = qbase(z0);
#pragma unroll
  for (int j = 0; j < 13; ++j) {
    double f;
    const int zz = zwrap(b + j, f);
    q[j] = vmul(f, P[(size_t)zz * plane + col]);
  }
  // value of plane zp (queue base bq) at this thread's column, from the queue
  // (uniform slot: 6 or 7 inside the box, others within 6 planes of its ends
  // in mode 0; registers only)
  auto centre = [&](int slot) {
    double2 v = q[6];
#pragma unroll
    for (int j = 0; j < 13; ++j)
      if (slot == j) v = q[j];
    return v;
  };
  // plane z0 of the tile into buffer 0
  {
    double2 hv[HU];
#pragma unroll
    for (int u = 0; u < HU; ++u) hv[u] = P[(size_t)z0 * plane + hsrc[u]];
    buf[0][cpos] = centre(z0 - b);
#pragma unroll
    for (int u = 0; u < HU; ++u)
      if (hoff[u] >= 0) buf[0][hoff[u]] = vmul(hsg[u], hv[u]);
  }
  // the halos of planes z0+1..z0+6 into the ring
  {
    double2 hp[6][HU];
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const size_t pj = (size_t)min(z0 + 1 + j, z1 - 1) * plane;
#pragma unroll
      for (int u = 0; u < HU; ++u) hp[j][u] = P[pj + hsrc[u]];
    }
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      double2* R = ring + ((z0 + 1 + j) % 6) * NHALO;
#pragma unroll
      for (int u = 0; u < HU; ++u)
        if (hoff[u] >= 0) R[threadIdx.x + u * NT] = vmul(hsg[u], hp[j][u]);
    }
  }
  // the epilogue's operands at the point, loaded one step ahead
  typename Out::Ops ops = o.load((size_t)z0 * plane + col, gx, gy, z0);
  // Per step: the halo of plane z+7 (for the ring) and the next window value
  // are loaded before the barrier; after the stencils the next plane is
  // staged into the other LDS buffer from the window and the ring; the
  // epilogue runs on operands loaded in the previous step; the next step's
  // operands are loaded, and the halo of plane z+7 goes into the ring last,
  // in the slot plane z+1 has just left.  Every load is issued
  // unconditionally (from a clamped, valid address when its value is not
  // needed), so each wait covers only loads issued well before it.
  for (int z = z0; z < z1; ++z) {
    const int zn = z + 1;
    const bool more = zn < z1;
    const int bn = more ? qbase(zn) : b;
    // (past the chunk: the chunk's last plane again, an L2 hit)
    const int zl = min(zn, z1 - 1);
    const size_t pn = (size_t)zl * plane;
    const size_t p7 = (size_t)min(z + 7, z1 - 1) * plane;
    double2 hv[HU];
#pragma unroll
    for (int u = 0; u < HU; ++u) hv[u] = P[p7 + hsrc[u]];
    double nqf;
    const int nqz = zwrap(bn + 12, nqf);
    const double2 nqr = P[(size_t)nqz * plane + col];
    __syncthreads();
    const double2* T = buf[(z - z0) & 1];
    const int k = z - b;
    const double2 cen = centre(k);
    const double2 dx = stencil_tile<double2>(cfx, T + (yl + 6) * TW, 1, gx, nx, x0);
    const double2 dy = stencil_tile<double2>(cfy, T + 6 + xl, TW, gy, ny, y0);
    double2 dz = make_double2(0.0, 0.0);
    if (k == 6) {
#pragma unroll
      for (int j = 0; j < 13; ++j) dz = vfma(cfz.c[j], q[j], dz);
    } else {
      const int r = (z < 6) ? z : z - nz + 12;
#pragma unroll
      for (int j = 0; j < 13; ++j) dz = vfma(cfz.rows[r][j], q[j], dz);
    }
    double2* R = ring + (zn % 6) * NHALO;
    if (more) {
      double2* Tn = buf[(zn - z0) & 1];
      Tn[cpos] = centre(zn - b);
#pragma unroll
      for (int u = 0; u < HU; ++u)
        if (hoff[u] >= 0) Tn[hoff[u]] = R[threadIdx.x + u * NT];
    }
    const double2 sxy = make_double2(dx.x + dy.x, dx.y + dy.y);
    o.apply((size_t)z * plane + col, gx, gy, z, make_double2(sxy.x + dz.x, sxy.y + dz.y), cen, ops, red);
    ops = o.load(pn + col, gx, gy, zl);
    // plane z+7 takes the slot of plane z+1 (z+7 = z+1 mod 6)
#pragma unroll
    for (int u = 0; u < HU; ++u)
      if (hoff[u] >= 0) R[threadIdx.x + u * NT] = vmul(hsg[u], hv[u]);
    if (bn != b) {
#pragma unroll
      for (int j = 0; j < 12; ++j) q[j] = q[j + 1];
      q[12] = vmul(nqf, nqr);
      b = bn;
    }
  }
  if constexpr (NR > 0) block_reduce_store<NR>(red, out.partials);
}

// geometry of the 2.5-D kernel: tile (TX, TY) = (32, 8) when the planes allow
// it, else (16, 16); 0 when neither divides the plane
struct Lap25Geom {
  int tx, ty, tz;
  size_t nblocks;
};
inline Lap25Geom lap25_geom(const Dims& d, int tz) {
  Lap25Geom g{0, 0, tz, 0};
  if (d.nx < 13 || d.ny < 13 || d.nz < 13 || tz < 1) return g;
  if (d.nx % 32 == 0 && d.ny % 8 == 0) {
    g.tx = 32;
    g.ty = 8;
  } else if (d.nx % 16 == 0 && d.ny % 16 == 0) {
    g.tx = 16;
    g.ty = 16;
  } else {
    return g;
  }
  g.nblocks = (size_t)(d.nx / g.tx) * (d.ny / g.ty) * ((d.nz + tz - 1) / tz);
  return g;
}

template <class Out>
inline void launch_lap25(hipStream_t s, const Dims& d, const StencilCoef* cf3, const double2* psi, const Out& o,
                         int tz) {
  const Lap25Geom g = lap25_geom(d, tz);
  if (g.tx == 32)
    hipLaunchKernelGGL((k_lap25<32, 8, Out>), dim3(g.nblocks), dim3(256), 0, s, psi, o, cf3, d, tz);
  else
    hipLaunchKernelGGL((k_lap25<16, 16, Out>), dim3(g.nblocks), dim3(256), 0, s, psi, o, cf3, d, tz);
}

// ---- common functors ---------------------------------------------------------
// Loads return the LDS staging type; stores receive the double accumulator and
// round to the storage type.
template <typename V>
struct InPlain {
  const V* a;
  __device__ __forceinline__ V load(size_t i) const { return a[i]; }
  __device__ __forceinline__ V loadx(size_t i) const { return a[i]; }
};
// (1 - denalf/den0s) * int_a   (term_alfa.f90:80-85)
template <typename R>
struct InFInt {
  const R *denalf, *ia;
  double den0s;
  __device__ __forceinline__ R load(size_t i) const { return (R)((1.0 - (double)denalf[i] / den0s) * ia[i]); }
};
// (f int_x, f int_y) for the fused x + y divergence pass
template <typename R>
struct InFInt2 {
  const R *denalf, *ix, *iy;
  double den0s;
  __device__ __forceinline__ R loadx(size_t i) const { return (R)((1.0 - (double)denalf[i] / den0s) * ix[i]); }
  __device__ __forceinline__ R loady(size_t i) const { return (R)((1.0 - (double)denalf[i] / den0s) * iy[i]); }
};
__device__ __forceinline__ double2 vadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double vadd(double a, double b) { return a + b; }
template <typename V>
__device__ __forceinline__ V vround(acc_t<V> a);
template <>
__device__ __forceinline__ double vround<double>(double a) { return a; }
template <>
__device__ __forceinline__ double2 vround<double2>(double2 a) { return a; }
template <>
__device__ __forceinline__ float vround<float>(double a) { return (float)a; }
template <>
__device__ __forceinline__ float2 vround<float2>(double2 a) { return make_float2((float)a.x, (float)a.y); }
// out = dx + dy (Laplacian / divergence partial sums)
template <typename V>
struct OutSumXY {
  static constexpr int NR = 0;
  V* out;
  double* partials;
  __device__ __forceinline__ void store(size_t i, int, int, int, acc_t<V> dx, acc_t<V> dy, double*) const {
    out[i] = vround<V>(vadd(dx, dy));
  }
};
// two separate outputs (derden)
template <typename R>
struct OutTwo {
  static constexpr int NR = 0;
  R *ox, *oy;
  double* partials;
  __device__ __forceinline__ void store(size_t i, int, int, int, double dx, double dy, double*) const {
    ox[i] = (R)dx;
    oy[i] = (R)dy;
  }
};
template <typename V>
struct OutStore {
  static constexpr int NR = 0;
  V* out;
  double* partials;
  __device__ __forceinline__ void store(size_t i, int, int, int, acc_t<V> v, double*) const {
    out[i] = vround<V>(v);
  }
};

// ---- 2.5-D divergence of f int_a with the mean-field assembly ---------------
// div = d_x(f int_x) + d_y(f int_y) + d_z(f int_z), f = 1 - denalf/den0s
// (term_alfa.f90:80-85), then ualphas = (conv/den0s + div) h2o2m4 alphas and
// pot4 = delj4 + dencg^2 (cp4/2 + cpp4/3 dencg) + corr [+ ext] + ualphas
// (term_alfa.f90:86-98, Dynamics_TDDFT/poten.f90:194-211).  Tiling as k_lap25:
// a workgroup owns TX x TY (x, y) columns over a chunk of at most tz planes;
// each thread keeps the 13-plane z window of f int_z (storage type R) in
// registers, and each plane of f int_x (with 6-column x halos) and f int_y
// (with 6-row y halos) is staged in LDS, double-buffered.  The operands are
// rounded to R where the separate passes stage them, dx + dy is rounded to R
// as the x+y pass stores it, and every sum keeps its order: the results equal
// those of k_pencil_xy2<InFInt2, OutSumXY> followed by k_pencil<InFInt,
// OutDivPot4>.  Requires nx % TX == 0, ny % TY == 0, TX, TY >= 7, n >= 13.
// The mean-field assembly at a point, shared by k_div25 and the z pass of the
// two-stream mean field so that both round identically (no contraction into
// fma is left to the compiler, which could fuse differently in the two
// kernels): ualphas = (conv/den0s + div) halpha and
// pot4 = delj4 + dencg^2 (cp4/2 + cpp4/3 dencg) + corr [+ ext] + ualphas.
struct AlphaPot4 {
  double ua, pot4;
};
__device__ __forceinline__ AlphaPot4 alpha_pot4(double conv, double den0s, double div, double halpha,
                                                double delj4, double cg, double corr, double a0, double a1,
                                                bool lext, double ext) {
#pragma clang fp contract(off)
  const double ua = (conv / den0s + div) * halpha;
  double p = delj4 + cg * cg * (a0 + a1 * cg) + corr;
  if (lext) p += ext;
  return AlphaPot4{ua, p + ua};
}

template <typename R>
struct Div25Args {
  const R *denalf, *ix, *iy, *iz;
  const R *conv, *delj4, *dencg, *corr, *ext;
  R *ualphas, *pot4;
  double den0s, halpha, a0, a1;
  int lext;
};

template <int TX, int TY, typename R>
__global__ void __launch_bounds__(TX * TY) k_div25(Div25Args<R> a, const StencilCoef* __restrict__ cf3, Dims d,
                                                  int tz) {
  static_assert(TX >= 7 && TY >= 7, "edge tiles must hold 13 rows");
  constexpr int NT = TX * TY, TW = TX + 12, TH = TY + 12;
  // LDS per buffer: the x operand (TY rows x TW, column 0 <-> x0 - 6), then
  // the y operand (TH rows x TX, row 0 <-> y0 - 6)
  constexpr int XS = TY * TW, BS = XS + TH * TX;
  // halo slots: the y operand's 6-row bands above and below the tile, then
  // the x operand's 6-column bands left and right of it
  constexpr int NHALO = 12 * TX + 12 * TY, HU = (NHALO + NT - 1) / NT;
  __shared__ R buf[2][BS];
  __shared__ StencilCoef cfs[3];
  static_assert(sizeof(StencilCoef) % 8 == 0, "copied as doubles");
  for (int e = threadIdx.x; e < (int)(3 * sizeof(StencilCoef) / 8); e += NT)
    reinterpret_cast<double*>(cfs)[e] = reinterpret_cast<const double*>(cf3)[e];
  __syncthreads();
  const StencilCoef& cfx = cfs[0];
  const StencilCoef& cfy = cfs[1];
  const StencilCoef& cfz = cfs[2];
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const int ntx = nx / TX, nty = ny / TY, nzc = (nz + tz - 1) / tz;
  // XCD-aware order (as k_lap25)
  const int ntiles = ntx * nty * nzc;
  int tile = blockIdx.x;
  if ((ntiles & 7) == 0) tile = (tile & 7) * (ntiles >> 3) + (tile >> 3);
  const int x0 = (tile % ntx) * TX, y0 = ((tile / ntx) % nty) * TY, zc = tile / (ntx * nty);
  const int z0 = zc * tz, z1 = min(z0 + tz, nz);
  const int xl = threadIdx.x % TX, yl = threadIdx.x / TX;
  const int gx = x0 + xl, gy = y0 + yl;
  const size_t plane = (size_t)nx * ny;
  const size_t col = (size_t)gy * nx + gx;
  const int xpos = yl * TW + xl + 6, ypos = XS + (yl + 6) * TX + xl;
  const double den0s = a.den0s;
  // f int at index i of one operand, rounded to R as the separate passes stage it
  auto fint = [&](const R* ia, size_t i) { return (R)((1.0 - (double)a.denalf[i] / den0s) * ia[i]); };
  const double sgx = (cfx.mode == 2) ? -1.0 : 1.0, sgy = (cfy.mode == 2) ? -1.0 : 1.0;
  const double sgz = (cfz.mode == 2) ? -1.0 : 1.0;
  int hoff[HU];
  size_t hsrc[HU];
  double hsg[HU];
  const R* hop[HU];
#pragma unroll
  for (int u = 0; u < HU; ++u) {
    const int k = min((int)threadIdx.x + u * NT, NHALO - 1);
    double sg = 1.0;
    int xx, yy, off;
    if (k < 12 * TX) {  // y operand: rows 0..5 and TY+6..TY+11
      const int r = (k < 6 * TX) ? k / TX : TY + 6 + (k - 6 * TX) / TX, c = k % TX;
      yy = y0 - 6 + r;
      xx = x0 + c;
      if (yy < 0 || yy >= ny) {
        if (cfy.mode != 0) {
          sg = sgy;
          yy += (yy < 0) ? ny : -ny;
        } else {
          yy = min(max(yy, 0), ny - 1);
        }
      }
      off = XS + r * TX + c;
      hop[u] = a.iy;
    } else {  // x operand: columns 0..5 and TX+6..TX+11
      const int k3 = k - 12 * TX, r = k3 / 12, cc = k3 % 12, c = (cc < 6) ? cc : TX + cc;
      yy = y0 + r;
      xx = x0 - 6 + c;
      if (xx < 0 || xx >= nx) {
        if (cfx.mode != 0) {
          sg = sgx;
          xx += (xx < 0) ? nx : -nx;
        } else {
          xx = min(max(xx, 0), nx - 1);
        }
      }
      off = r * TW + c;
      hop[u] = a.ix;
    }
    hoff[u] = ((int)threadIdx.x + u * NT < NHALO) ? off : -1;
    hsrc[u] = (size_t)yy * nx + xx;
    hsg[u] = sg;
  }
  const bool clampz = cfz.mode == 0;
  auto qbase = [&](int z) { return clampz ? min(max(z - 6, 0), nz - 13) : z - 6; };
  auto zwrap = [&](int u, double& f) {
    f = 1.0;
    if (u < 0) {
      f = sgz;
      return u + nz;
    }
    if (u >= nz) {
      f = sgz;
      return u - nz;
    }
    return u;
  };
  R q[13];
  int b = qbase(z0);
#pragma unroll
  for (int j = 0; j < 13; ++j) {
    double f;
    const int zz = zwrap(b + j, f);
    q[j] = vmul(f, fint(a.iz, (size_t)zz * plane + col));
  }
  // plane z0 of the tile into buffer 0
  {
    const size_t p0 = (size_t)z0 * plane;
    R hv[HU];
#pragma unroll
    for (int u = 0; u < HU; ++u) hv[u] = fint(hop[u], p0 + hsrc[u]);
    buf[0][xpos] = fint(a.ix, p0 + col);
    buf[0][ypos] = fint(a.iy, p0 + col);
#pragma unroll
    for (int u = 0; u < HU; ++u)
      if (hoff[u] >= 0) buf[0][hoff[u]] = vmul(hsg[u], hv[u]);
  }
  // the assembly's operands at the point, loaded one step ahead (ext is read
  // from conv's address when there is no external term)
  const R* ext = a.lext ? a.ext : a.conv;
  struct Ops {
    double cv, dj, cg, co, ex;
  };
  auto load_ops = [&](size_t i) {
    return Ops{(double)a.conv[i], (double)a.delj4[i], (double)a.dencg[i], (double)a.corr[i], (double)ext[i]};
  };
  Ops ops = load_ops((size_t)z0 * plane + col);
  for (int z = z0; z < z1; ++z) {
    const int zn = z + 1;
    const bool more = zn < z1;
    const int bn = more ? qbase(zn) : b;
    const size_t pn = (size_t)min(zn, nz - 1) * plane;
    R hv[HU];
#pragma unroll
    for (int u = 0; u < HU; ++u) hv[u] = fint(hop[u], pn + hsrc[u]);
    const R cxn = fint(a.ix, pn + col), cyn = fint(a.iy, pn + col);
    double nqf;
    const int nqz = zwrap(bn + 12, nqf);
    const R nqr = fint(a.iz, (size_t)nqz * plane + col);
    __syncthreads();
    const R* T = buf[(z - z0) & 1];
    const double dx = stencil_tile<R>(cfx, T + yl * TW, 1, gx, nx, x0);
    const double dy = stencil_tile<R>(cfy, T + XS + xl, TX, gy, ny, y0);
    const int k = z - b;
    double dz = 0.0;
    if (k == 6) {
#pragma unroll
      for (int j = 0; j < 13; ++j) dz = vfma(cfz.c[j], q[j], dz);
    } else {
      const int r = (z < 6) ? z : z - nz + 12;
#pragma unroll
      for (int j = 0; j < 13; ++j) dz = vfma(cfz.rows[r][j], q[j], dz);
    }
    if (more) {
      R* Tn = buf[(zn - z0) & 1];
      Tn[xpos] = cxn;
      Tn[ypos] = cyn;
#pragma unroll
      for (int u = 0; u < HU; ++u)
        if (hoff[u] >= 0) Tn[hoff[u]] = vmul(hsg[u], hv[u]);
    }
    // OutSumXY's rounding of dx + dy, then OutDivPot4's arithmetic
    const double sxy = (double)vround<R>(dx + dy);
    const size_t i = (size_t)z * plane + col;
    const AlphaPot4 r = alpha_pot4(ops.cv, den0s, sxy + dz, a.halpha, ops.dj, ops.cg, ops.co, a.a0, a.a1,
                                   a.lext != 0, ops.ex);
    a.ualphas[i] = (R)r.ua;
    a.pot4[i] = (R)r.pot4;
    ops = load_ops(pn + col);
    if (bn != b) {
#pragma unroll
      for (int j = 0; j < 12; ++j) q[j] = q[j + 1];
      q[12] = vmul(nqf, nqr);
      b = bn;
    }
  }
}

template <typename R>
inline void launch_div25(hipStream_t s, const Dims& d, const StencilCoef* cf3, const Div25Args<R>& a, int tz) {
  const Lap25Geom g = lap25_geom(d, tz);
  if (g.tx == 32)
    hipLaunchKernelGGL((k_div25<32, 8, R>), dim3(g.nblocks), dim3(256), 0, s, a, cf3, d, tz);
  else
    hipLaunchKernelGGL((k_div25<16, 16, R>), dim3(g.nblocks), dim3(256), 0, s, a, cf3, d, tz);
}

}  // namespace hedft
