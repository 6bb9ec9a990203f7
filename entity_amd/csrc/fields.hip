// fields.hip -- Minkowski Yee stencils, current source term, digital filter,
// periodic ghost fill and periodic current fold.
//
// One thread per active cell, x fastest (64-wide rows), so that every
// wavefront reads and writes contiguous 256-byte segments of each component
// plane.  These kernels are HBM-bound.
#include "internal.hpp"

namespace ntt {
namespace {

constexpr int BX = 64, BY = 4;

struct CellIdx {
  int i, j, k;
  bool ok;
};

__device__ __forceinline__ CellIdx active_cell(const FieldView& F) {
  CellIdx c;
  const int a = blockIdx.x * BX + threadIdx.x;
  const int b = blockIdx.y * BY + threadIdx.y;
  const int g = blockIdx.z;
  c.ok = (a < F.n1) && (b < F.n2);
  c.i  = a + F.ngh;
  c.j  = b + F.ngh;
  c.k  = (F.N3 > 1) ? g + F.ngh : 0;
  return c;
}

dim3 active_grid(const FieldView& F) {
  return dim3((F.n1 + BX - 1) / BX, (F.n2 + BY - 1) / BY, F.N3 > 1 ? F.n3 : 1);
}

// ---- Faraday: src/kernels/faraday_mink.hpp:71-166 ---------------------------
struct FD {
  float c1, c2;
  float dx, dy, dz, bxy, byx, bxz, bzx, byz, bzy;
};

template <bool EXT>
__global__ __launch_bounds__(BX* BY) void k_faraday_2d(FieldView F, FD f) {
  const CellIdx c = active_cell(F);
  if (!c.ok) return;
  const int i1 = c.i, i2 = c.j;
  auto E = [&](int a, int b, int comp) -> float { return F(a, b, 0, comp); };
  if constexpr (!EXT) {
    // standard Yee: alpha = 1, delta = beta = 0
    F(i1, i2, 0, 3) += f.c1 * (-(E(i1, i2 + 1, 2) - E(i1, i2, 2)));
    F(i1, i2, 0, 4) += f.c1 * (+(E(i1 + 1, i2, 2) - E(i1, i2, 2)));
    F(i1, i2, 0, 5) += f.c2 * (+(E(i1, i2 + 1, 0) - E(i1, i2, 0)) - (E(i1 + 1, i2, 1) - E(i1, i2, 1)));
  } else {
    const float alphax = ONE - TWO * f.bxy - THREE * f.dx;
    const float alphay = ONE - TWO * f.byx - THREE * f.dy;
    F(i1, i2, 0, 3) += f.c1 * (-alphay * (E(i1, i2 + 1, 2) - E(i1, i2, 2)) -
                               f.dy * (E(i1, i2 + 2, 2) - E(i1, i2 - 1, 2)) -
                               f.byx * (E(i1 + 1, i2 + 1, 2) - E(i1 + 1, i2, 2)) -
                               f.byx * (E(i1 - 1, i2 + 1, 2) - E(i1 - 1, i2, 2)));
    F(i1, i2, 0, 4) += f.c1 * (+alphax * (E(i1 + 1, i2, 2) - E(i1, i2, 2)) +
                               f.dx * (E(i1 + 2, i2, 2) - E(i1 - 1, i2, 2)) +
                               f.bxy * (E(i1 + 1, i2 + 1, 2) - E(i1, i2 + 1, 2)) +
                               f.bxy * (E(i1 + 1, i2 - 1, 2) - E(i1, i2 - 1, 2)));
    F(i1, i2, 0, 5) += f.c2 * (+alphay * (E(i1, i2 + 1, 0) - E(i1, i2, 0)) +
                               f.dy * (E(i1, i2 + 2, 0) - E(i1, i2 - 1, 0)) +
                               f.byx * (E(i1 + 1, i2 + 1, 0) - E(i1 + 1, i2, 0)) +
                               f.byx * (E(i1 - 1, i2 + 1, 0) - E(i1 - 1, i2, 0)) -
                               alphax * (E(i1 + 1, i2, 1) - E(i1, i2, 1)) -
                               f.dx * (E(i1 + 2, i2, 1) - E(i1 - 1, i2, 1)) -
                               f.bxy * (E(i1 + 1, i2 + 1, 1) - E(i1, i2 + 1, 1)) -
                               f.bxy * (E(i1 + 1, i2 - 1, 1) - E(i1, i2 - 1, 1)));
  }
}

template <bool EXT>
__global__ __launch_bounds__(BX* BY) void k_faraday_3d(FieldView F, FD f) {
  const CellIdx c = active_cell(F);
  if (!c.ok) return;
  const int i1 = c.i, i2 = c.j, i3 = c.k;
  auto E = [&](int a, int b, int g, int comp) -> float { return F(a, b, g, comp); };
  if constexpr (!EXT) {
    F(i1, i2, i3, 3) += f.c1 * (+(E(i1, i2, i3 + 1, 1) - E(i1, i2, i3, 1)) -
                                (E(i1, i2 + 1, i3, 2) - E(i1, i2, i3, 2)));
    F(i1, i2, i3, 4) += f.c1 * (+(E(i1 + 1, i2, i3, 2) - E(i1, i2, i3, 2)) -
                                (E(i1, i2, i3 + 1, 0) - E(i1, i2, i3, 0)));
    F(i1, i2, i3, 5) += f.c1 * (+(E(i1, i2 + 1, i3, 0) - E(i1, i2, i3, 0)) -
                                (E(i1 + 1, i2, i3, 1) - E(i1, i2, i3, 1)));
  } else {
    const float alphax = ONE - TWO * f.bxy - TWO * f.bxz - THREE * f.dx;
    const float alphay = ONE - TWO * f.byx - TWO * f.byz - THREE * f.dy;
    const float alphaz = ONE - TWO * f.bzx - TWO * f.bzy - THREE * f.dz;
    F(i1, i2, i3, 3) += f.c1 * (
      +alphaz * (E(i1, i2, i3 + 1, 1) - E(i1, i2, i3, 1))
      + f.dz * (E(i1, i2, i3 + 2, 1) - E(i1, i2, i3 - 1, 1))
      + f.bzx * (E(i1 + 1, i2, i3 + 1, 1) - E(i1 + 1, i2, i3, 1))
      + f.bzx * (E(i1 - 1, i2, i3 + 1, 1) - E(i1 - 1, i2, i3, 1))
      + f.bzy * (E(i1, i2 + 1, i3 + 1, 1) - E(i1, i2 + 1, i3, 1))
      + f.bzy * (E(i1, i2 - 1, i3 + 1, 1) - E(i1, i2 - 1, i3, 1))
      - alphay * (E(i1, i2 + 1, i3, 2) - E(i1, i2, i3, 2))
      - f.dy * (E(i1, i2 + 2, i3, 2) - E(i1, i2 - 1, i3, 2))
      - f.byx * (E(i1 + 1, i2 + 1, i3, 2) - E(i1 + 1, i2, i3, 2))
      - f.byx * (E(i1 - 1, i2 + 1, i3, 2) - E(i1 - 1, i2, i3, 2))
      - f.byz * (E(i1, i2 + 1, i3 + 1, 2) - E(i1, i2, i3 + 1, 2))
      - f.byz * (E(i1, i2 + 1, i3 - 1, 2) - E(i1, i2, i3 - 1, 2)));
    F(i1, i2, i3, 4) += f.c1 * (
      +alphax * (E(i1 + 1, i2, i3, 2) - E(i1, i2, i3, 2))
      + f.dx * (E(i1 + 2, i2, i3, 2) - E(i1 - 1, i2, i3, 2))
      + f.bxy * (E(i1 + 1, i2 + 1, i3, 2) - E(i1, i2 + 1, i3, 2))
      + f.bxy * (E(i1 + 1, i2 - 1, i3, 2) - E(i1, i2 - 1, i3, 2))
      + f.bxz * (E(i1 + 1, i2, i3 + 1, 2) - E(i1, i2, i3 + 1, 2))
      + f.bxz * (E(i1 + 1, i2, i3 - 1, 2) - E(i1, i2, i3 - 1, 2))
      - alphaz * (E(i1, i2, i3 + 1, 0) - E(i1, i2, i3, 0))
      - f.dz * (E(i1, i2, i3 + 2, 0) - E(i1, i2, i3 - 1, 0))
      - f.bzx * (E(i1 + 1, i2, i3 + 1, 0) - E(i1 + 1, i2, i3, 0))
      - f.bzx * (E(i1 - 1, i2, i3 + 1, 0) - E(i1 - 1, i2, i3, 0))
      - f.bzy * (E(i1, i2 + 1, i3 + 1, 0) - E(i1, i2 + 1, i3, 0))
      - f.bzy * (E(i1, i2 - 1, i3 + 1, 0) - E(i1, i2 - 1, i3, 0)));
    F(i1, i2, i3, 5) += f.c1 * (
      +alphay * (E(i1, i2 + 1, i3, 0) - E(i1, i2, i3, 0))
      + f.dy * (E(i1, i2 + 2, i3, 0) - E(i1, i2 - 1, i3, 0))
      + f.byx * (E(i1 + 1, i2 + 1, i3, 0) - E(i1 + 1, i2, i3, 0))
      + f.byx * (E(i1 - 1, i2 + 1, i3, 0) - E(i1 - 1, i2, i3, 0))
      + f.byz * (E(i1, i2 + 1, i3 + 1, 0) - E(i1, i2, i3 + 1, 0))
      + f.byz * (E(i1, i2 + 1, i3 - 1, 0) - E(i1, i2, i3 - 1, 0))
      - alphax * (E(i1 + 1, i2, i3, 1) - E(i1, i2, i3, 1))
      - f.dx * (E(i1 + 2, i2, i3, 1) - E(i1 - 1, i2, i3, 1))
      - f.bxy * (E(i1 + 1, i2 + 1, i3, 1) - E(i1, i2 + 1, i3, 1))
      - f.bxy * (E(i1 + 1, i2 - 1, i3, 1) - E(i1, i2 - 1, i3, 1))
      - f.bxz * (E(i1 + 1, i2, i3 + 1, 1) - E(i1, i2, i3 + 1, 1))
      - f.bxz * (E(i1 + 1, i2, i3 - 1, 1) - E(i1, i2, i3 - 1, 1)));
  }
}

// ---- Ampere: src/kernels/ampere_mink.hpp:57-86 ------------------------------
__global__ __launch_bounds__(BX* BY) void k_ampere_2d(FieldView F, float c1, float c2) {
  const CellIdx c = active_cell(F);
  if (!c.ok) return;
  const int i1 = c.i, i2 = c.j;
  F(i1, i2, 0, 0) += c1 * (F(i1, i2, 0, 5) - F(i1, i2 - 1, 0, 5));
  F(i1, i2, 0, 1) += c1 * (F(i1 - 1, i2, 0, 5) - F(i1, i2, 0, 5));
  F(i1, i2, 0, 2) += c2 * (F(i1, i2 - 1, 0, 3) - F(i1, i2, 0, 3) + F(i1, i2, 0, 4) - F(i1 - 1, i2, 0, 4));
}

__global__ __launch_bounds__(BX* BY) void k_ampere_3d(FieldView F, float c1) {
  const CellIdx c = active_cell(F);
  if (!c.ok) return;
  const int i1 = c.i, i2 = c.j, i3 = c.k;
  F(i1, i2, i3, 0) += c1 * (F(i1, i2, i3 - 1, 4) - F(i1, i2, i3, 4) + F(i1, i2, i3, 5) -
                            F(i1, i2 - 1, i3, 5));
  F(i1, i2, i3, 1) += c1 * (F(i1 - 1, i2, i3, 5) - F(i1, i2, i3, 5) + F(i1, i2, i3, 3) -
                            F(i1, i2, i3 - 1, 3));
  F(i1, i2, i3, 2) += c1 * (F(i1, i2 - 1, i3, 3) - F(i1, i2, i3, 3) + F(i1, i2, i3, 4) -
                            F(i1 - 1, i2, i3, 4));
}

// ---- E += coeff * J; J /= ppc0: src/kernels/ampere_mink.hpp:156-210 ---------
__global__ __launch_bounds__(BX* BY) void k_currents_ampere(FieldView E, FieldView J, float coeff,
                                                            float ppc0) {
  const CellIdx c = active_cell(E);
  if (!c.ok) return;
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    const float j = J(c.i, c.j, c.k, comp);
    E(c.i, c.j, c.k, comp) += j * coeff;
    J(c.i, c.j, c.k, comp) = j / ppc0;
  }
}

// ---- digital filter, periodic interior: digital_filter.hpp:182-194, 349-378
__global__ __launch_bounds__(BX* BY) void k_filter_2d(FieldView A, FieldView B) {
  const CellIdx c = active_cell(A);
  if (!c.ok) return;
  const int i1 = c.i, i2 = c.j;
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    auto b = [&](int x, int y) -> float { return B(x, y, 0, comp); };
    A(i1, i2, 0, comp) = INV_4 * b(i1, i2) +
                         INV_8 * (b(i1 - 1, i2) + b(i1 + 1, i2) + b(i1, i2 - 1) + b(i1, i2 + 1)) +
                         INV_16 * (b(i1 - 1, i2 - 1) + b(i1 + 1, i2 + 1) + b(i1 - 1, i2 + 1) +
                                   b(i1 + 1, i2 - 1));
  }
}

// 3D: the 1/32 group is the reference's list as written (it names
// (i,j,k-1) and (i,j,k+1) where the edge pair (i,j-1,k+1), (i,j+1,k-1) would
// complete the binomial stencil).
__global__ __launch_bounds__(BX* BY) void k_filter_3d(FieldView A, FieldView B) {
  const CellIdx c = active_cell(A);
  if (!c.ok) return;
  const int i = c.i, j = c.j, k = c.k;
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    auto b = [&](int x, int y, int z) -> float { return B(x, y, z, comp); };
    A(i, j, k, comp) =
      INV_8 * b(i, j, k) +
      INV_16 * (b(i - 1, j, k) + b(i + 1, j, k) + b(i, j - 1, k) + b(i, j + 1, k) + b(i, j, k - 1) +
                b(i, j, k + 1)) +
      INV_32 * (b(i - 1, j - 1, k) + b(i + 1, j + 1, k) + b(i - 1, j + 1, k) + b(i + 1, j - 1, k) +
                b(i, j - 1, k - 1) + b(i, j + 1, k + 1) + b(i, j, k - 1) + b(i, j, k + 1) +
                b(i - 1, j, k - 1) + b(i + 1, j, k + 1) + b(i - 1, j, k + 1) + b(i + 1, j, k - 1)) +
      INV_64 * (b(i - 1, j - 1, k - 1) + b(i + 1, j + 1, k + 1) + b(i - 1, j + 1, k + 1) +
                b(i + 1, j - 1, k - 1) + b(i - 1, j - 1, k + 1) + b(i + 1, j + 1, k - 1) +
                b(i - 1, j + 1, k - 1) + b(i + 1, j - 1, k + 1));
  }
}

// ---- periodic ghost fill: metadomain_comm.cpp:204-367 -----------------------
__device__ __forceinline__ int wrap_src(int i, int G, int n) {
  return i < G ? i + n : (i >= G + n ? i - n : i);
}

__global__ __launch_bounds__(BX* BY) void k_ghost_fill(FieldView F, int c0, int c1) {
  const int i = blockIdx.x * BX + threadIdx.x;
  const int j = blockIdx.y * BY + threadIdx.y;
  const int k = blockIdx.z;
  if (i >= F.N1 || j >= F.N2) return;
  const int G  = F.ngh;
  const int si = wrap_src(i, G, F.n1);
  const int sj = wrap_src(j, G, F.n2);
  const int sk = (F.N3 > 1) ? wrap_src(k, G, F.n3) : k;
  if (si == i && sj == j && sk == k) return;
  for (int comp = c0; comp < c1; ++comp) F(i, j, k, comp) = F(si, sj, sk, comp);
}

// ---- periodic J fold: metadomain_comm.cpp:408-562 ---------------------------
// For an active cell, every direction d whose receive slab holds it adds the
// cell's periodic image; images are summed in Directions<D>::all order
// (src/global/arch/directions.h:180-233) and the sum is added to the cell.
__global__ __launch_bounds__(BX* BY) void k_fold(FieldView J) {
  const CellIdx c = active_cell(J);
  if (!c.ok) return;
  const int G = J.ngh;
  const int D = (J.N3 > 1) ? 3 : 2;
  const int n[3] = { J.n1, J.n2, J.n3 };
  const int x[3] = { c.i, c.j, c.k };
  int lo[3], hi[3];  // allowed direction components per axis
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (d >= D) { lo[d] = 0; hi[d] = 0; continue; }
    lo[d] = (x[d] >= n[d]) ? -1 : 0;          // receives from dir -1 : i in [n, n+G)
    hi[d] = (x[d] < 2 * G) ? 1 : 0;           // receives from dir +1 : i in [G, 2G)
  }
  if (lo[0] == 0 && hi[0] == 0 && lo[1] == 0 && hi[1] == 0 && lo[2] == 0 && hi[2] == 0) return;
  for (int comp = 0; comp < 3; ++comp) {
    float acc = ZERO;
    bool  any = false;
    for (int a = lo[0]; a <= hi[0]; ++a)
      for (int b = lo[1]; b <= hi[1]; ++b)
        for (int g = lo[2]; g <= hi[2]; ++g) {
          if (a == 0 && b == 0 && g == 0) continue;
          // direction (a,b,g): source = x + dir * n  (dir +1 -> ghost above)
          acc = any ? acc + J(x[0] + a * n[0], x[1] + b * n[1], x[2] + g * n[2], comp)
                    : ZERO + J(x[0] + a * n[0], x[1] + b * n[1], x[2] + g * n[2], comp);
          any = true;
        }
    J(x[0], x[1], x[2], comp) += acc;
  }
}

// ---- periodic-box variants used by the engine step ---------------------------
// Neighbours are read at their periodic image inside the active region, so the
// kernels do not depend on ghost cells; every thread writes its active cell
// and all of that cell's ghost copies, so the array leaves the kernel with
// consistent periodic ghosts (what the reference obtains with
// CommunicateFields afterwards).  The arithmetic is that of the kernels above.
struct PIdx {
  // active-image index of x + off (x ghost-extended active index)
  int G, n;
  __device__ __forceinline__ int at(int x, int off) const {
    int y = x - G + off;
    y += (y < 0) ? n : 0;
    y -= (y >= n) ? n : 0;
    return y + G;
  }
};

// writes v at (i, j, k) and at every ghost copy of that active cell
__device__ __forceinline__ void store_with_ghosts(const FieldView& F, int i, int j, int k, int c,
                                                  float v) {
  const int G = F.ngh;
  const int xi[2] = { i, (i - G < G) ? i + F.n1 : ((i - G >= F.n1 - G) ? i - F.n1 : -1) };
  const int xj[2] = { j, (j - G < G) ? j + F.n2 : ((j - G >= F.n2 - G) ? j - F.n2 : -1) };
  int       xk[2] = { k, -1 };
  if (F.N3 > 1) xk[1] = (k - G < G) ? k + F.n3 : ((k - G >= F.n3 - G) ? k - F.n3 : -1);
  for (int a = 0; a < 2; ++a) {
    if (xi[a] < 0) continue;
    for (int b = 0; b < 2; ++b) {
      if (xj[b] < 0) continue;
      for (int g = 0; g < 2; ++g) {
        if (xk[g] < 0) continue;
        F(xi[a], xj[b], xk[g], c) = v;
      }
    }
  }
}

__device__ __forceinline__ void faraday_periodic_cell(const FieldView& F, int i, int j, int k,
                                                      float c1, float c2) {
  const PIdx X{ F.ngh, F.n1 }, Y{ F.ngh, F.n2 }, Z{ F.ngh, F.n3 };
  const int ip = X.at(i, 1), jp = Y.at(j, 1);
  if (F.N3 == 1) {
    const float b1 = F(i, j, 0, 3) + c1 * (-(F(i, jp, 0, 2) - F(i, j, 0, 2)));
    const float b2 = F(i, j, 0, 4) + c1 * (+(F(ip, j, 0, 2) - F(i, j, 0, 2)));
    const float b3 = F(i, j, 0, 5) + c2 * (+(F(i, jp, 0, 0) - F(i, j, 0, 0)) - (F(ip, j, 0, 1) - F(i, j, 0, 1)));
    store_with_ghosts(F, i, j, 0, 3, b1);
    store_with_ghosts(F, i, j, 0, 4, b2);
    store_with_ghosts(F, i, j, 0, 5, b3);
  } else {
    const int kp = Z.at(k, 1);
    const float b1 = F(i, j, k, 3) + c1 * (+(F(i, j, kp, 1) - F(i, j, k, 1)) - (F(i, jp, k, 2) - F(i, j, k, 2)));
    const float b2 = F(i, j, k, 4) + c1 * (+(F(ip, j, k, 2) - F(i, j, k, 2)) - (F(i, j, kp, 0) - F(i, j, k, 0)));
    const float b3 = F(i, j, k, 5) + c1 * (+(F(i, jp, k, 0) - F(i, j, k, 0)) - (F(ip, j, k, 1) - F(i, j, k, 1)));
    store_with_ghosts(F, i, j, k, 3, b1);
    store_with_ghosts(F, i, j, k, 4, b2);
    store_with_ghosts(F, i, j, k, 5, b3);
  }
}

__global__ __launch_bounds__(BX* BY) void k_faraday_periodic(FieldView F, float c1, float c2) {
  const CellIdx c = active_cell(F);
  if (!c.ok) return;
  faraday_periodic_cell(F, c.i, c.j, c.k, c1, c2);
}

// Ampere (E += coeff * curl B) followed by the current term (E += J * coeff,
// J /= ppc0) in one pass; with_j = false skips the current term.
__device__ __forceinline__ void ampere_currents_periodic_cell(const FieldView& F, const FieldView& J,
                                                              int i, int j, int k, float c1,
                                                              float c2, float coeff, float ppc0,
                                                              int with_j) {
  const PIdx X{ F.ngh, F.n1 }, Y{ F.ngh, F.n2 }, Z{ F.ngh, F.n3 };
  const int im = X.at(i, -1), jm = Y.at(j, -1);
  float e[3];
  if (F.N3 == 1) {
    e[0] = F(i, j, 0, 0) + c1 * (F(i, j, 0, 5) - F(i, jm, 0, 5));
    e[1] = F(i, j, 0, 1) + c1 * (F(im, j, 0, 5) - F(i, j, 0, 5));
    e[2] = F(i, j, 0, 2) + c2 * (F(i, jm, 0, 3) - F(i, j, 0, 3) + F(i, j, 0, 4) - F(im, j, 0, 4));
  } else {
    const int km = Z.at(k, -1);
    e[0] = F(i, j, k, 0) + c1 * (F(i, j, km, 4) - F(i, j, k, 4) + F(i, j, k, 5) - F(i, jm, k, 5));
    e[1] = F(i, j, k, 1) + c1 * (F(im, j, k, 5) - F(i, j, k, 5) + F(i, j, k, 3) - F(i, j, km, 3));
    e[2] = F(i, j, k, 2) + c1 * (F(i, jm, k, 3) - F(i, j, k, 3) + F(i, j, k, 4) - F(im, j, k, 4));
  }
  for (int comp = 0; comp < 3; ++comp) {
    if (with_j) {
      const float jv = J(i, j, k, comp);
      e[comp] += jv * coeff;
      store_with_ghosts(J, i, j, k, comp, jv / ppc0);
    }
    store_with_ghosts(F, i, j, k, comp, e[comp]);
  }
}

__global__ __launch_bounds__(BX* BY) void k_ampere_currents_periodic(FieldView F, FieldView J,
                                                                     float c1, float c2,
                                                                     float coeff, float ppc0,
                                                                     int with_j) {
  const CellIdx c = active_cell(F);
  if (!c.ok) return;
  ampere_currents_periodic_cell(F, J, c.i, c.j, c.k, c1, c2, coeff, ppc0, with_j);
}

// The 3D binomial stencil as written in digital_filter.hpp:349-378 (the 1/32
// group lists (0,0,-1) and (0,0,+1) where the edge pair (0,-1,+1), (0,+1,-1)
// would be), from the 27 source values b[dz + 1][dy + 1][dx + 1], in the
// reference's operation order.  Every filter kernel evaluates it here, so the
// compiler's contraction of the sum is the same in all of them.
__device__ __forceinline__ float filter3d_value(const float (&b)[3][3][3]) {
  return INV_8 * b[1][1][1] +
         INV_16 * (b[1][1][0] + b[1][1][2] + b[1][0][1] + b[1][2][1] + b[0][1][1] + b[2][1][1]) +
         INV_32 * (b[1][0][0] + b[1][2][2] + b[1][2][0] + b[1][0][2] + b[0][0][1] + b[2][2][1] +
                   b[0][1][1] + b[2][1][1] + b[0][1][0] + b[2][1][2] + b[2][1][0] + b[0][1][2]) +
         INV_64 * (b[0][0][0] + b[2][2][2] + b[2][2][0] + b[0][0][2] + b[2][0][0] + b[0][2][2] +
                   b[0][2][0] + b[2][0][2]);
}

// one filter pass dst = K * src (periodic reads), 2D and 3D
__device__ __forceinline__ void filter_periodic_cell(const FieldView& Dst, const FieldView& Src,
                                                     int i, int j, int k) {
  const PIdx X{ Src.ngh, Src.n1 }, Y{ Src.ngh, Src.n2 }, Z{ Src.ngh, Src.n3 };
  const int im = X.at(i, -1), ip = X.at(i, 1), jm = Y.at(j, -1), jp = Y.at(j, 1);
  for (int comp = 0; comp < 3; ++comp) {
    float v;
    if (Dst.N3 == 1) {
      auto b = [&](int x, int y) -> float { return Src(x, y, 0, comp); };
      v = INV_4 * b(i, j) + INV_8 * (b(im, j) + b(ip, j) + b(i, jm) + b(i, jp)) +
          INV_16 * (b(im, jm) + b(ip, jp) + b(im, jp) + b(ip, jm));
    } else {
      const int km = Z.at(k, -1), kp = Z.at(k, 1);
      const int xs[3] = { im, i, ip }, ys[3] = { jm, j, jp }, zs[3] = { km, k, kp };
      float b[3][3][3];
#pragma unroll
      for (int z = 0; z < 3; ++z)
#pragma unroll
        for (int y = 0; y < 3; ++y)
#pragma unroll
          for (int x = 0; x < 3; ++x) b[z][y][x] = Src(xs[x], ys[y], zs[z], comp);
      v = filter3d_value(b);
    }
    store_with_ghosts(Dst, i, j, k, comp, v);
  }
}

__global__ __launch_bounds__(BX* BY) void k_filter_periodic(FieldView Dst, FieldView Src) {
  const CellIdx c = active_cell(Dst);
  if (!c.ok) return;
  filter_periodic_cell(Dst, Src, c.i, c.j, c.k);
}

// The 3D pass of k_filter_periodic, tiled for gfx950: a workgroup owns a
// 64 x 4 column of cells and marches it through KC planes.  The three source
// planes a plane's stencil reads (each with a one-cell halo, 66 x 6 x 3
// components) sit in an LDS ring, so every source value is read from HBM once
// per workgroup instead of 27 times through the caches, and the 27 reads of
// each output come from LDS with immediate offsets.  Neighbours are read at
// their ghost-extended positions: the source's periodic ghosts are filled
// (the current from the fixed-point accumulator and every filter pass write
// them).  The expression and its operation order are those of
// filter_periodic_cell.
constexpr int FTX = 64, FTY = 4, FKC = 16;
constexpr int FHX = FTX + 2, FHY = FTY + 2, FPL = FHX * FHY;

// Each thread moves FLD fixed elements of every source plane (component,
// row, column: offsets computed once); the next plane's loads are issued into
// registers before the current plane's stencil is computed, so their latency
// overlaps it, and are written to the ring after.
constexpr int FLD = (3 * FPL + FTX * FTY - 1) / (FTX * FTY);

__global__ __launch_bounds__(FTX* FTY) void k_filter_periodic_3d(FieldView Dst, FieldView Src) {
  __shared__ float ring[3][3 * FPL];
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * FTX + tx;
  const int G = Src.ngh;
  const int i0 = blockIdx.x * FTX, j0 = blockIdx.y * FTY, k0 = blockIdx.z * FKC;
  const int kend = min(k0 + FKC, Src.n3);
  // this thread's elements of a plane: offset in the source (within a plane
  // of ghost-extended index 0; -1: outside the array, loaded as zero)
  int64_t goff[FLD];
#pragma unroll
  for (int u = 0; u < FLD; ++u) {
    const int e = tid + u * FTX * FTY;
    goff[u]     = -1;
    if (e < 3 * FPL) {
      const int c = e / FPL, r = e - c * FPL;
      const int hy = r / FHX, hx = r - hy * FHX;
      const int gi = i0 + hx - 1 + G, gj = j0 + hy - 1 + G;
      if (gi < Src.N1 && gj < Src.N2) goff[u] = Src.idx(gi, gj, 0, c);
    }
  }
  const int64_t pstride = static_cast<int64_t>(Src.N1) * Src.N2;
  float pre[FLD];
  // source plane kk (active index, may be -1 or n3: a ghost plane) into pre
  auto fetch = [&](int kk) {
    const float* base = Src.d + static_cast<int64_t>(kk + G) * pstride;
#pragma unroll
    for (int u = 0; u < FLD; ++u) pre[u] = (goff[u] >= 0) ? base[goff[u]] : ZERO;
  };
  auto commit = [&](int slot) {
#pragma unroll
    for (int u = 0; u < FLD; ++u) {
      const int e = tid + u * FTX * FTY;
      if (e < 3 * FPL) ring[slot][e] = pre[u];
    }
  };
  fetch(k0 - 1);
  commit((k0 + 2) % 3);
  fetch(k0);
  commit(k0 % 3);
  fetch(k0 + 1);
  const int i = i0 + tx, j = j0 + ty;
  const bool active = i < Src.n1 && j < Src.n2;
  const int  o = (ty + 1) * FHX + (tx + 1);
  for (int k = k0; k < kend; ++k) {
    commit((k + 1) % 3);
    __syncthreads();
    if (k + 2 <= kend) fetch(k + 2);
    if (active) {
      const int sm = (k + 2) % 3, s0 = k % 3, sp = (k + 1) % 3;
#pragma unroll
      for (int comp = 0; comp < 3; ++comp) {
        const float* pl[3] = { ring[sm] + comp * FPL, ring[s0] + comp * FPL,
                               ring[sp] + comp * FPL };
        float b[3][3][3];
#pragma unroll
        for (int z = 0; z < 3; ++z)
#pragma unroll
          for (int y = 0; y < 3; ++y)
#pragma unroll
            for (int x = 0; x < 3; ++x) b[z][y][x] = pl[z][o + (y - 1) * FHX + (x - 1)];
        store_with_ghosts(Dst, i + G, j + G, k + G, comp, filter3d_value(b));
      }
    }
    __syncthreads();
  }
}

__global__ void k_fill_zero(float* __restrict__ d, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t k = i; k < n; k += stride) d[k] = ZERO;
}

__global__ void k_fill_zero_u64(unsigned long long* __restrict__ d, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t k = i; k < n; k += stride) d[k] = 0ull;
}

// float current from the 64-bit fixed-point accumulator (units of 1/scale,
// a power of two): exact in double below 2^53, then one rounding to float
__device__ __forceinline__ void j_from_fixed_cell(const FieldView& J,
                                                  const unsigned long long* __restrict__ acc,
                                                  double inv_scale, int i, int j, int k) {
  for (int comp = 0; comp < 3; ++comp) {
    const long long q = static_cast<long long>(acc[J.idx(i, j, k, comp)]);
    const float v = static_cast<float>(static_cast<double>(q) * inv_scale);
    store_with_ghosts(J, i, j, k, comp, v);
  }
}

__global__ __launch_bounds__(BX* BY) void k_j_from_fixed(FieldView J,
                                                         const unsigned long long* __restrict__ acc,
                                                         double inv_scale) {
  const CellIdx c = active_cell(J);
  if (!c.ok) return;
  j_from_fixed_cell(J, acc, inv_scale, c.i, c.j, c.k);
}

__global__ void k_copy(float* __restrict__ dst, const float* __restrict__ src, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t k = i; k < n; k += stride) dst[k] = src[k];
}

// ---- batched launches over the boxes of an ensemble ---------------------------
// Segment s owns blocks [segs[s].blk0, segs[s + 1].blk0) of a 1D grid; a block
// finds its segment by binary search and runs the same cell function as the
// single-box kernel, so every box gets the arithmetic of its own launch.
template <class Seg>
__device__ __forceinline__ int find_seg(const Seg* __restrict__ segs, int nseg, int64_t b) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (segs[mid].blk0 <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int OP>
__global__ __launch_bounds__(BX* BY) void k_fields_batched(const FieldSeg* __restrict__ segs,
                                                           int nseg) {
  const int64_t b = blockIdx.x;
  const int     s = find_seg(segs, nseg, b);
  const FieldSeg g = segs[s];
  const int64_t lb = b - g.blk0;
  const int64_t per_z = static_cast<int64_t>(g.gx) * g.gy;
  const int bz = static_cast<int>(lb / per_z);
  const int rr = static_cast<int>(lb - bz * per_z);
  const int by = rr / g.gx, bx = rr - by * g.gx;
  const int a = bx * BX + threadIdx.x, bb = by * BY + threadIdx.y;
  const FieldView& F = g.F;
  if (a >= F.n1 || bb >= F.n2) return;
  const int i = a + F.ngh, j = bb + F.ngh, k = (F.N3 > 1) ? bz + F.ngh : 0;
  if constexpr (OP == SEG_FARADAY) faraday_periodic_cell(F, i, j, k, g.c1, g.c2);
  else if constexpr (OP == SEG_AMPERE)
    ampere_currents_periodic_cell(F, g.J, i, j, k, g.c1, g.c2, g.coeff, g.ppc0, g.with_j);
  else if constexpr (OP == SEG_FILTER) filter_periodic_cell(F, g.J, i, j, k);
  else j_from_fixed_cell(F, g.acc, g.inv_scale, i, j, k);
}

constexpr int kFlatBlock = 256, kFlatPerThread = 8;

template <int OP>
__global__ __launch_bounds__(kFlatBlock) void k_flat_batched(const FlatSeg* __restrict__ segs,
                                                             int nseg) {
  const int64_t b = blockIdx.x;
  const int     s = find_seg(segs, nseg, b);
  const FlatSeg g = segs[s];
  const int64_t base = (b - g.blk0) * (kFlatBlock * kFlatPerThread) + threadIdx.x;
#pragma unroll
  for (int r = 0; r < kFlatPerThread; ++r) {
    const int64_t e = base + static_cast<int64_t>(r) * kFlatBlock;
    if (e >= g.n) break;
    if constexpr (OP == FLAT_ZERO_U64) static_cast<unsigned long long*>(g.dst)[e] = 0ull;
    else static_cast<float*>(g.dst)[e] = static_cast<const float*>(g.src)[e];
  }
}

ntt_status launch_ok() { return check_launch(); }

}  // namespace
}  // namespace ntt

using namespace ntt;

extern "C" ntt_status ntt_faraday_mink(ntt_field_t* em, float coeff1, float coeff2,
                                       const ntt_fd_coeffs_t* fd, void* stream) {
  FieldView F;
  ntt_status s = make_field_view(em, 6, &F);
  if (s != NTT_OK) return s;
  FD f{ coeff1, coeff2, 0, 0, 0, 0, 0, 0, 0, 0, 0 };
  bool ext = false;
  if (fd != nullptr) {
    f.dx = fd->delta_x; f.dy = fd->delta_y; f.dz = fd->delta_z;
    f.bxy = fd->beta_xy; f.byx = fd->beta_yx; f.bxz = fd->beta_xz;
    f.bzx = fd->beta_zx; f.byz = fd->beta_yz; f.bzy = fd->beta_zy;
    ext = f.dx != 0 || f.dy != 0 || f.dz != 0 || f.bxy != 0 || f.byx != 0 || f.bxz != 0 ||
          f.bzx != 0 || f.byz != 0 || f.bzy != 0;
  }
  if (ext && F.ngh < 2) return NTT_ERR_ARG;
  const dim3 grid = active_grid(F), block(BX, BY);
  hipStream_t st = as_stream(stream);
  if (em->dim == 2) {
    if (ext) k_faraday_2d<true><<<grid, block, 0, st>>>(F, f);
    else k_faraday_2d<false><<<grid, block, 0, st>>>(F, f);
  } else {
    if (ext) k_faraday_3d<true><<<grid, block, 0, st>>>(F, f);
    else k_faraday_3d<false><<<grid, block, 0, st>>>(F, f);
  }
  return launch_ok();
}

extern "C" ntt_status ntt_ampere_mink(ntt_field_t* em, float coeff1, float coeff2, void* stream) {
  FieldView F;
  ntt_status s = make_field_view(em, 6, &F);
  if (s != NTT_OK) return s;
  const dim3 grid = active_grid(F), block(BX, BY);
  if (em->dim == 2) k_ampere_2d<<<grid, block, 0, as_stream(stream)>>>(F, coeff1, coeff2);
  else k_ampere_3d<<<grid, block, 0, as_stream(stream)>>>(F, coeff1);
  return launch_ok();
}

static bool same_shape(const FieldView& a, const FieldView& b) {
  return a.n1 == b.n1 && a.n2 == b.n2 && a.n3 == b.n3 && a.ngh == b.ngh && a.N3 == b.N3;
}

extern "C" ntt_status ntt_currents_ampere_mink(ntt_field_t* em, ntt_field_t* cur, float coeff,
                                               float ppc0, void* stream) {
  FieldView E, J;
  ntt_status s = make_field_view(em, 6, &E);
  if (s != NTT_OK) return s;
  s = make_field_view(cur, 3, &J);
  if (s != NTT_OK) return s;
  if (!same_shape(E, J)) return NTT_ERR_ARG;
  k_currents_ampere<<<active_grid(E), dim3(BX, BY), 0, as_stream(stream)>>>(E, J, coeff, ppc0);
  return launch_ok();
}

extern "C" ntt_status ntt_digital_filter(ntt_field_t* array, const ntt_field_t* buffer,
                                         const ntt_flds_bc_t* bc, void* stream) {
  FieldView A, B;
  ntt_status s = make_field_view(array, 3, &A);
  if (s != NTT_OK) return s;
  s = make_field_view(buffer, 3, &B);
  if (s != NTT_OK) return s;
  if (!same_shape(A, B) || A.d == B.d) return NTT_ERR_ARG;
  if (bc != nullptr) {
    // conductor and axis special cases (digital_filter.hpp:129-180) are not built
    for (int d = 0; d < array->dim; ++d)
      for (int e = 0; e < 2; ++e)
        if (bc->bc[d][e] != NTT_FLDS_BC_PERIODIC) return NTT_ERR_UNSUPPORTED;
  }
  if (array->dim == 2) k_filter_2d<<<active_grid(A), dim3(BX, BY), 0, as_stream(stream)>>>(A, B);
  else k_filter_3d<<<active_grid(A), dim3(BX, BY), 0, as_stream(stream)>>>(A, B);
  return launch_ok();
}

extern "C" ntt_status ntt_communicate_fields(ntt_field_t* fld, int32_t comp_begin, int32_t comp_end,
                                             void* stream) {
  FieldView F;
  ntt_status s = make_field_view(fld, 0, &F);
  if (s != NTT_OK) return s;
  if (comp_begin < 0 || comp_end > F.nc || comp_begin >= comp_end) return NTT_ERR_ARG;
  if (F.n1 < F.ngh || F.n2 < F.ngh || (F.N3 > 1 && F.n3 < F.ngh)) return NTT_ERR_ARG;
  const dim3 grid((F.N1 + BX - 1) / BX, (F.N2 + BY - 1) / BY, F.N3);
  k_ghost_fill<<<grid, dim3(BX, BY), 0, as_stream(stream)>>>(F, comp_begin, comp_end);
  return launch_ok();
}

extern "C" ntt_status ntt_synchronize_currents(ntt_field_t* cur, ntt_field_t* buff, void* stream) {
  (void)buff;
  FieldView J;
  ntt_status s = make_field_view(cur, 3, &J);
  if (s != NTT_OK) return s;
  if (J.n1 < 2 * J.ngh || J.n2 < 2 * J.ngh || (J.N3 > 1 && J.n3 < 2 * J.ngh)) return NTT_ERR_ARG;
  k_fold<<<active_grid(J), dim3(BX, BY), 0, as_stream(stream)>>>(J);
  return launch_ok();
}

// ---- engine-internal launchers ------------------------------------------------
namespace ntt {

unsigned flat_grid(int64_t n) {
  int64_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return static_cast<unsigned>(g);
}

void launch_faraday_periodic(const FieldView& F, float c1, float c2, hipStream_t st) {
  k_faraday_periodic<<<active_grid(F), dim3(BX, BY), 0, st>>>(F, c1, c2);
}

void launch_ampere_currents_periodic(const FieldView& F, const FieldView& J, float c1, float c2,
                                     float coeff, float ppc0, bool with_j, hipStream_t st) {
  k_ampere_currents_periodic<<<active_grid(F), dim3(BX, BY), 0, st>>>(F, J, c1, c2, coeff, ppc0,
                                                                     with_j ? 1 : 0);
}

void launch_filter_periodic(const FieldView& dst, const FieldView& src, hipStream_t st) {
  if (dst.N3 > 1) {
    const dim3 grid((dst.n1 + FTX - 1) / FTX, (dst.n2 + FTY - 1) / FTY, (dst.n3 + FKC - 1) / FKC);
    k_filter_periodic_3d<<<grid, dim3(FTX, FTY), 0, st>>>(dst, src);
    return;
  }
  k_filter_periodic<<<active_grid(dst), dim3(BX, BY), 0, st>>>(dst, src);
}

void launch_ghost_fill(const FieldView& F, int c0, int c1, hipStream_t st) {
  const dim3 grid((F.N1 + BX - 1) / BX, (F.N2 + BY - 1) / BY, F.N3);
  k_ghost_fill<<<grid, dim3(BX, BY), 0, st>>>(F, c0, c1);
}

void launch_fold(const FieldView& J, hipStream_t st) {
  k_fold<<<active_grid(J), dim3(BX, BY), 0, st>>>(J);
}

void launch_zero(float* d, int64_t n, hipStream_t st) {
  k_fill_zero<<<flat_grid(n), 256, 0, st>>>(d, n);
}

void launch_copy(float* dst, const float* src, int64_t n, hipStream_t st) {
  k_copy<<<flat_grid(n), 256, 0, st>>>(dst, src, n);
}

void launch_zero_u64(unsigned long long* d, int64_t n, hipStream_t st) {
  k_fill_zero_u64<<<flat_grid(n), 256, 0, st>>>(d, n);
}

void launch_j_from_fixed(const FieldView& J, const unsigned long long* acc, double inv_scale,
                         hipStream_t st) {
  k_j_from_fixed<<<active_grid(J), dim3(BX, BY), 0, st>>>(J, acc, inv_scale);
}

int64_t field_seg_blocks(const FieldView& F, int32_t* gx, int32_t* gy) {
  const dim3 g = active_grid(F);
  *gx = static_cast<int32_t>(g.x);
  *gy = static_cast<int32_t>(g.y);
  return static_cast<int64_t>(g.x) * g.y * g.z;
}

void launch_fields_batched(int op, const FieldSeg* segs, int nseg, int64_t nblocks,
                           hipStream_t st) {
  if (nseg == 0 || nblocks == 0) return;
  const dim3 grid(static_cast<unsigned>(nblocks)), block(BX, BY);
  if (op == SEG_FARADAY) k_fields_batched<SEG_FARADAY><<<grid, block, 0, st>>>(segs, nseg);
  else if (op == SEG_AMPERE) k_fields_batched<SEG_AMPERE><<<grid, block, 0, st>>>(segs, nseg);
  else if (op == SEG_FILTER) k_fields_batched<SEG_FILTER><<<grid, block, 0, st>>>(segs, nseg);
  else k_fields_batched<SEG_JFIX><<<grid, block, 0, st>>>(segs, nseg);
}

int64_t flat_seg_blocks(int64_t n) {
  const int64_t per = static_cast<int64_t>(kFlatBlock) * kFlatPerThread;
  return (n + per - 1) / per;
}

void launch_flat_batched(int op, const FlatSeg* segs, int nseg, int64_t nblocks, hipStream_t st) {
  if (nseg == 0 || nblocks == 0) return;
  if (op == FLAT_ZERO_U64)
    k_flat_batched<FLAT_ZERO_U64><<<static_cast<unsigned>(nblocks), kFlatBlock, 0, st>>>(segs, nseg);
  else
    k_flat_batched<FLAT_COPY_F32><<<static_cast<unsigned>(nblocks), kFlatBlock, 0, st>>>(segs, nseg);
}

}  // namespace ntt
